/*
 * camera.h - camera state of the renderer, mirroring the reference's
 * CameraParameters (src/camera/cameraParameters.h:18-129), Camera
 * (src/camera/camera.h:8-35) and the non-UI part of CameraController
 * (src/camera/cameraController.h:7-26).  Mouse and key handling of the
 * orbit/fly controllers belongs to the UI and is not part of this layer; the
 * orbit and fly motions themselves are available as explicit calls.
 *
 * Coordinate frame (as in the reference): forward = normalize(target - eye),
 * right = normalize((0,1,0) x forward), up = forward x right.
 */
#pragma once

#include "vt_math_types.h"

namespace voxeltoy {

class CameraController;

class CameraParameters {
public:
    CameraParameters();

    enum CameraLensModel { CLM_PINHOLE, CLM_THIN_LENS, CLM_ORTHOGRAPHIC };

    const V3f& eye() const { return m_eye; }
    const V3f& target() const { return m_target; }
    float distanceToTarget() const;

    void getBasis(V3f& forwardUnitVector, V3f& rightUnitVector, V3f& upUnitVector) const;
    V3f forwardUnitVector() const;
    V3f rightUnitVector() const;
    V3f upUnitVector() const;

    /* polar angle from +Y and azimuth from +X toward +Z of the view direction */
    float rotationTheta() const;
    float rotationPhi() const;

    float fovY() const { return m_fovY; }
    /* film height / (2 tan(fovY/2)) */
    float focalLength() const;
    float nearDistance() const { return m_near; }
    float farDistance() const { return m_far; }
    float focalDistance() const { return m_focalDistance; }
    V2f filmSize() const { return m_filmSize; }
    float lensRadius() const { return m_lensRadius; }
    CameraLensModel lensModel() const { return m_lensModel; }

    /* 35 mm film: 36 mm wide */
    static constexpr float FILM_SIZE_35MM = 36.0f;

private:
    friend class CameraController;
    friend class Camera;

    void lookAt(const V3f& target);
    void setDistanceFromTarget(float distance);
    void orbitAroundTarget(float theta, float phi);
    void orbitAroundEye(float theta, float phi);
    void setEyeTarget(const V3f& eye, const V3f& target);
    void setFovY(float fov);
    void setNearDistance(float d);
    void setFarDistance(float d);
    void setFocalLength(float length);
    void setFocalDistance(float distance);
    void setFilmSize(float filmW, float filmH);
    void setLensRadius(float radius);
    void setFStop(float fstop);
    void setLensModel(CameraLensModel model);

    V3f m_target, m_eye;
    float m_fovY = 0.0f;
    float m_near, m_far;
    float m_focalDistance;
    float m_lensRadius;
    V2f m_filmSize;
    CameraLensModel m_lensModel = CLM_PINHOLE;
};

/* The controller's programmatic operations (cameraController.cpp:4-24),
 * plus the orbit and fly motions the UI controllers drive, and the input
 * handling of the two controllers the reference switches between
 * (orbitCameraController.cpp:15-46, flyCameraController.cpp:15-78).
 * Mouse motion is in image-normalised units; buttons and keys use Qt's
 * codes (right button 0x2, middle button 0x4, letter keys = upper-case
 * ASCII). */
class CameraController {
public:
    enum Mode { ORBIT, FLY };
    explicit CameraController(CameraParameters* p) : m_parameters(p) {}
    /* ORBIT: right button orbits the eye around the target, middle button
     * dollies by 1.05x.  FLY: right button turns the view around the eye,
     * middle button moves eye and target along the up vector (100 units
     * per unit of dy).  Returns whether the camera changed. */
    bool onMouseMove(float dx, float dy, int buttons);
    /* FLY: W/S move forward/back, D/A right/left by 100 units.  ORBIT
     * takes no keys. */
    bool onKeyPress(int key);
    Mode mode() const { return m_mode; }
    void setMode(Mode m) { m_mode = m; }
    void lookAt(const V3f& target);
    void setDistanceFromTarget(float distance);
    /* centre the bounds and back off so their diagonal fills the vertical fov */
    void focusOnBounds(const Box3f& bounds);
    /* eye on the sphere around the target at polar angle theta (from +Y) and
     * azimuth phi, distance unchanged */
    void orbit(float theta, float phi);
    /* target on the sphere around the eye (fly-camera look) */
    void look(float theta, float phi);
    void setEyeTarget(const V3f& eye, const V3f& target);

private:
    CameraParameters* m_parameters;
    Mode m_mode = ORBIT;
};

class Camera {
public:
    Camera() : m_controller(&m_parameters) {}
    Camera(const Camera& o) : m_parameters(o.m_parameters), m_controller(&m_parameters)
    {
        m_controller.setMode(o.m_controller.mode());
    }
    Camera& operator=(const Camera& o)
    {
        m_parameters = o.m_parameters;
        m_controller.setMode(o.m_controller.mode());
        return *this;
    }

    const CameraParameters& parameters() const { return m_parameters; }
    CameraController& controller() { return m_controller; }

    void setLensModel(CameraParameters::CameraLensModel m) { m_parameters.setLensModel(m); }
    void setFocalLength(float length) { m_parameters.setFocalLength(length); }
    void setFocalDistance(float distance) { m_parameters.setFocalDistance(distance); }
    void setFilmSize(float filmW, float filmH) { m_parameters.setFilmSize(filmW, filmH); }
    void setLensRadius(float radius) { m_parameters.setLensRadius(radius); }
    void setFStop(float fstop) { m_parameters.setFStop(fstop); }
    void setFovY(float fov) { m_parameters.setFovY(fov); }
    void setNearDistance(float d) { m_parameters.setNearDistance(d); }
    void setFarDistance(float d) { m_parameters.setFarDistance(d); }

    /* Camera::CameraControllerMode (src/camera/camera.h:22-27): which
     * controller interprets mouse and key input; a new camera orbits. */
    enum CameraControllerMode { CCM_ORBIT, CCM_FLY };
    void setCameraController(CameraControllerMode mode)
    {
        m_controller.setMode(mode == CCM_FLY ? CameraController::FLY : CameraController::ORBIT);
    }
    CameraControllerMode cameraControllerMode() const
    {
        return m_controller.mode() == CameraController::FLY ? CCM_FLY : CCM_ORBIT;
    }

private:
    CameraParameters m_parameters;
    CameraController m_controller;
};

}  // namespace voxeltoy
